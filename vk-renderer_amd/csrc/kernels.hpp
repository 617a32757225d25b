// kernels.hpp - device data layout shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vkr.h"
#include "bvh.hpp"

namespace vkr {

constexpr uint32_t kMiss = 0xFFFFFFFFu;

// Everything a kernel needs of one uploaded scene (device pointers).
struct DevScene {
    const BvhNode* nodes;
    const TriRecord* tris;          // leaf order
    const float4* sph;              // leaf order: centre, r
    const uint32_t* sph_key;        // leaf order: sphere scan index
    const vkr_triangle* tris_orig;  // scan order (std430, as uploaded)
    const vkr_sphere* spheres;      // scan order
    const vkr_plane* planes;        // scan order
    const vkr_material* tri_material;
    const vkr_material* materials;
    const uint32_t* emitters;       // scan indices of spheres with emissive != 0
    uint32_t n_tris, n_spheres, n_planes, n_emitters;
    uint32_t n_materials, flags, has_bvh, stack_needed;
    uint32_t emitters_bounded;      // every emitter: |emissive| < 2^20, 2^-20 <= r <= 2^20
    float sph_c[3], sph_r;          // every sphere centre lies within sph_r of sph_c
};

// One frame of work (a single trace call, or one job of a batch).
struct JobDesc {
    DevScene sc;
    float cam_pos[3];
    float aspect;
    float cam_dir[3];
    float seed;
    float cam_right[3];
    float eps;
    float cam_up[3];
    uint32_t flags;               // VKR_TRACE_*
    float light_pos[3];           // FrameData.light_pos (read by the Whitted kernel only)
    uint32_t kernel;              // VKR_KERNEL_*
    uint32_t W, H, spp, max_depth;
    uint32_t tiles_x, n_tiles, pool_base, pad;
    uint8_t* rgba8;
    float* radiance;
    uint32_t* prim_id;
    float* hit_t;
    uint32_t* ray_count;
    uint16_t* decisions;
    uint2* adj_samples;           // W*H*spp, pixel-major: (bits of the random-stream position at the
                                  // sample's start, clamp-saturation bits) for the adjoint; NULL: none
};

struct DevCounters {
    unsigned long long rays, shadow, traced, nodes, tris, spheres, planes, overflow;
    unsigned long long wave_nodes, wave_leaves, wave_steps;
};

// Default pool refill threshold: a wave claims new pixels once at least this
// many of its lanes are idle (8-16 measured best at 1080p; 32 idles lanes,
// 1 refills too often).
constexpr uint32_t kDefaultChunk = 8u;

// Most work items one launch's 32-bit pool cursor hands out.  A resident
// wave can add up to 64 to the cursor after the last item has gone (its one
// claim that finds the pool empty), so the cursor stays below 2^32 while the
// resident lanes number under 2^24.
constexpr uint64_t kMaxPoolItems = 0xFF000000ull;

// One-pixel-per-lane launches cover at most this many lanes each, which
// bounds their ray records (2^21 x kLaneBytes = 2.7 GB).
constexpr uint32_t kMaxPlainLanes = 1u << 21;

// Ray records of the pooled path kernel, in global memory: one block of
// kWaveBytes per wave (64 lanes), laid out field by field so that the lanes
// of a wave reading or writing one field touch contiguous addresses.  Field
// f of lane l is at f + 16 l; per-slot fields hold slot k of lane l at
// f + 16 (64 k + l) (salts: f + 4 (64 k + l)).
//   kWPark   4 x float4: path state parked while the wave traces
//            (ro, dj.x) (rd, dj.y) (P, dj.z) (N, 0), at kWPark + 1024 i
//   kWHead   this step's ray origin xyz; w holds the bits of (job << 1) | 1
//            for shadow rays, (job << 1) for a closest-hit ray (job: index
//            in the batch, 0 for a single frame)
//   kWSlotD  per slot: the ray's (d xyz, t_max); slot k holds emitter
//            ei0 + k of a next-event window, or slot 0 the path segment
//   kWSlotC  per slot: the next-event contribution (xyz, 0)
//   kWHit    3 x float4: the diffuse hit published for the wave's
//            next-event evaluation: (N xyz, u), (albedo xyz, roughness),
//            (metalness, v, bits of ei0, bits of the kNee* flags)
//   kWSalt   per slot: the random stream position before emitter ei0 + k's
//            sample
//   kWPix    the sum of the pixel's finished samples (xyz)
constexpr uint32_t kRaySlots = 32;
constexpr uint32_t kWPark = 0u;
constexpr uint32_t kWHead = kWPark + 4u * 1024u;
constexpr uint32_t kWSlotD = kWHead + 1024u;
constexpr uint32_t kWSlotC = kWSlotD + kRaySlots * 1024u;
constexpr uint32_t kWHit = kWSlotC + kRaySlots * 1024u;
constexpr uint32_t kWSalt = kWHit + 3u * 1024u;
constexpr uint32_t kWPix = kWSalt + kRaySlots * 256u;
constexpr uint32_t kWaveBytes = kWPix + 1024u;
constexpr uint32_t kLaneBytes = kWaveBytes / 64u;   // what one lane of a launch needs
static_assert(kWaveBytes % 64u == 0u, "a wave's records split evenly over its lanes");

struct PathLaunch {
    JobDesc job;                  // single-frame launches
    const JobDesc* jobs;          // batched launches
    const uint32_t* job_start;    // n_jobs + 1 pool offsets (batched)
    uint32_t n_jobs;
    uint32_t total;               // pool indices (64 per 8x8 tile)
    uint32_t persistent;          // 0: one pixel per lane; 1: refill from pool
    uint32_t base;                // one pixel per lane: pool index of the launch's lane 0
    uint32_t chunk;               // refill when at least this many lanes idle
    uint32_t* pool;               // persistent pool cursor (device)
    DevCounters* counters;
    char* lanes;                  // ray records: one kWaveBytes block per wave of the grid
};

// Adjoint launch (path replay), see adjoint.hip.
struct AdjointLaunch {
    JobDesc job;
    const float* d_radiance;      // W*H*4
    float* d_params;              // 7 * n_spheres + 1
    const float* params;          // 7 * n_spheres + 1 (current values)
    DevCounters* counters;
    uint32_t* pool;               // persistent pool cursors: [0] pass 1, [1] pass 2
    uint32_t total;               // pool indices (64 per 8x8 tile; pass 2: per tile and sample)
    uint2* samples;               // W*H*spp: (bits of the random-stream position at the sample's
                                  // start, the pass-1 clamp-saturation bits), pixel-major
};

// One instantiation of the path kernel.  pool: a wave traces up to
// kRaySlots rays per lane and step from a shared pool (pays off with several
// emitters); else one ray per lane and step.  Results are identical.
struct PathVariant {
    uint32_t block, stack;
    bool count, batch, pool;
};

hipError_t launch_path(const PathLaunch& a, const PathVariant& v, uint32_t grid, hipStream_t s);
hipError_t launch_whitted(const JobDesc& j, DevCounters* c, hipStream_t s);
hipError_t launch_filter(const uint8_t* cur, const uint8_t* prev, uint32_t sw, uint32_t sh, uint8_t* out,
                         uint32_t dw, uint32_t dh, hipStream_t s);
// pass1 false: a.samples already holds the first pass's records of this frame.
hipError_t launch_adjoint(const AdjointLaunch& a, uint32_t stack, bool pass1, hipStream_t s);
hipError_t launch_debug(uint32_t op, const float* in, float* out, uint32_t n, hipStream_t s);
hipError_t launch_zero_counters(DevCounters* c, uint32_t* pool, hipStream_t s);
// Writes the per-frame fields of a job in a device job table (camera,
// aspect, seed, light position), in stream order.
hipError_t launch_set_frame(JobDesc* job, const vkr_frame_data& fd, hipStream_t s);
// Occupancy-derived persistent grid size for (block, stack).
uint32_t persistent_grid(const PathVariant& v, int device);

// The static scratch of vkr_trace_frame_static: counters at offset 0, the
// pool cursor at kStaticPoolOff, per-lane ray records from kStaticLanesOff.
constexpr size_t kStaticPoolOff = 256;
constexpr size_t kStaticLanesOff = 512;
constexpr size_t kStaticScratchBytes = kStaticLanesOff + (size_t)VKR_STATIC_SCRATCH_LANE_BYTES;
static_assert(sizeof(DevCounters) <= kStaticPoolOff, "counters fit below the pool cursor");
// Device address of this device's static scratch (looked up once per device
// with that device current, then cached).
hipError_t static_scratch(int device, char** out);
// Largest stack depth with a compiled kernel variant (0 when none fits).
uint32_t stack_variant(uint32_t needed);

// Device build of a triangle subtree (lbvh.hip): internal nodes go to
// d_nodes[node_off .. node_off + n - 2] (its root at node_off), one record
// per triangle to d_recs (Morton order; a leaf is a run of at most max_leaf
// records).  Returns, after a stream synchronisation, the subtree's height
// (its deepest leaf, the root's children at 1), its live node and leaf
// counts, and the exact bounds of the padded triangle boxes.  A tree higher
// than height_cap is rejected after its topology: only the height and counts
// are returned and the nodes' boxes and the records are not written.
struct LbvhRoot {
    float lo[3], hi[3];
    uint32_t height, nodes, leaves;
};
// Device scratch a build over n triangles needs (0 when n < 2); the caller
// passes at least that much to lbvh_build and may reuse it across builds.
size_t lbvh_scratch_bytes(uint32_t n);
hipError_t lbvh_build(const vkr_triangle* d_tris, uint32_t n, uint32_t node_off, uint32_t max_leaf,
                      uint32_t height_cap, void* scratch, size_t scratch_bytes, BvhNode* d_nodes, TriRecord* d_recs,
                      LbvhRoot* root, hipStream_t s);

}  // namespace vkr
