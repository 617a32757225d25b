// scene_file.cpp - scenes on disk (host only, no device is touched).
//
// The reference has no scene file: its scene is shader constants plus one
// triangle copied into the SceneData SSBO (Source/GraphicsDevice.cpp:796-822).
// A file here holds the arrays of a vkr_scene_desc in exactly those std430
// layouts behind a 64-byte header, so a scene written once (a synthetic
// workload, an exported mesh) can be replayed by any caller.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/vkr.h"
#include "ctx.hpp"

using namespace vkr;

namespace {

constexpr char kMagic[8] = {'V', 'K', 'R', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kVersion = 1;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t flags;
    uint32_t n_tris, n_spheres, n_planes, n_materials;
    uint32_t has_tri_material;
    uint32_t max_leaf_size, max_depth;
    float traversal_cost;
    uint32_t reserved[4];
};
static_assert(sizeof(FileHeader) == 64, "scene file header is 64 bytes");

uint64_t payload_bytes(const FileHeader& h)
{
    return (h.has_tri_material ? sizeof(vkr_material) : 0u) + (uint64_t)h.n_tris * sizeof(vkr_triangle) +
           (uint64_t)h.n_spheres * sizeof(vkr_sphere) + (uint64_t)h.n_planes * sizeof(vkr_plane) +
           (uint64_t)h.n_materials * sizeof(vkr_material);
}

struct File {
    FILE* f = nullptr;
    explicit File(FILE* p) : f(p) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

bool put(FILE* f, const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

}  // namespace

extern "C" {

vkr_error vkr_scene_save(const char* path, const vkr_scene_desc* d)
{
    if (!path || !d) return fail(VKR_INVALID_ARGUMENT, "NULL path or scene");
    if ((d->n_tris && !d->tris) || (d->n_spheres && !d->spheres) || (d->n_planes && !d->planes) ||
        (d->n_materials && !d->materials))
        return fail(VKR_INVALID_ARGUMENT, "a scene array is NULL but its count is not 0");
    FileHeader h{};
    std::memcpy(h.magic, kMagic, sizeof kMagic);
    h.version = kVersion;
    h.flags = d->flags;
    h.n_tris = d->n_tris;
    h.n_spheres = d->n_spheres;
    h.n_planes = d->n_planes;
    h.n_materials = d->n_materials;
    h.has_tri_material = d->tri_material ? 1u : 0u;
    h.max_leaf_size = d->max_leaf_size;
    h.max_depth = d->max_depth;
    h.traversal_cost = d->traversal_cost;
    File f(std::fopen(path, "wb"));
    if (!f.f) return fail(VKR_INVALID_ARGUMENT, std::string("cannot open ") + path + " for writing");
    const bool ok = put(f.f, &h, sizeof h) &&
                    (!d->tri_material || put(f.f, d->tri_material, sizeof(vkr_material))) &&
                    put(f.f, d->tris, (size_t)d->n_tris * sizeof(vkr_triangle)) &&
                    put(f.f, d->spheres, (size_t)d->n_spheres * sizeof(vkr_sphere)) &&
                    put(f.f, d->planes, (size_t)d->n_planes * sizeof(vkr_plane)) &&
                    put(f.f, d->materials, (size_t)d->n_materials * sizeof(vkr_material));
    if (!ok || std::fflush(f.f) != 0) return fail(VKR_UNKNOWN, std::string("write failed: ") + path);
    return VKR_SUCCESS;
}

vkr_error vkr_scene_load(const char* path, vkr_scene_desc* out)
{
    if (!path || !out) return fail(VKR_INVALID_ARGUMENT, "NULL path or output");
    *out = vkr_scene_desc{};
    File f(std::fopen(path, "rb"));
    if (!f.f) return fail(VKR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    FileHeader h{};
    if (!get(f.f, &h, sizeof h) || std::memcmp(h.magic, kMagic, sizeof kMagic) != 0)
        return fail(VKR_INVALID_ARGUMENT, std::string(path) + " is not a vkr scene file");
    if (h.version != kVersion) return fail(VKR_UNSUPPORTED, "unknown scene file version");
    if (h.has_tri_material > 1u) return fail(VKR_INVALID_ARGUMENT, "corrupt scene file header");
    if (std::fseek(f.f, 0, SEEK_END) != 0) return fail(VKR_UNKNOWN, "cannot size the scene file");
    const long end = std::ftell(f.f);
    if (end < 0 || (uint64_t)end != sizeof h + payload_bytes(h))
        return fail(VKR_INVALID_ARGUMENT, "scene file size does not match its header (truncated or corrupt)");
    if (std::fseek(f.f, (long)sizeof h, SEEK_SET) != 0) return fail(VKR_UNKNOWN, "cannot read the scene file");

    auto take = [&](size_t bytes, void** dst) -> bool {
        if (bytes == 0) return true;
        void* p = std::malloc(bytes);
        if (!p) return false;
        *dst = p;
        return get(f.f, p, bytes);
    };
    void *tm = nullptr, *tris = nullptr, *sph = nullptr, *pln = nullptr, *mats = nullptr;
    const bool ok = take(h.has_tri_material ? sizeof(vkr_material) : 0u, &tm) &&
                    take((size_t)h.n_tris * sizeof(vkr_triangle), &tris) &&
                    take((size_t)h.n_spheres * sizeof(vkr_sphere), &sph) &&
                    take((size_t)h.n_planes * sizeof(vkr_plane), &pln) &&
                    take((size_t)h.n_materials * sizeof(vkr_material), &mats);
    out->tri_material = static_cast<const vkr_material*>(tm);
    out->tris = static_cast<const vkr_triangle*>(tris);
    out->spheres = static_cast<const vkr_sphere*>(sph);
    out->planes = static_cast<const vkr_plane*>(pln);
    out->materials = static_cast<const vkr_material*>(mats);
    if (!ok) {
        vkr_scene_free(out);
        return fail(VKR_OUT_OF_MEMORY, "reading the scene file failed");
    }
    out->n_tris = h.n_tris;
    out->n_spheres = h.n_spheres;
    out->n_planes = h.n_planes;
    out->n_materials = h.n_materials;
    out->flags = h.flags;
    out->max_leaf_size = h.max_leaf_size;
    out->max_depth = h.max_depth;
    out->traversal_cost = h.traversal_cost;
    return VKR_SUCCESS;
}

void vkr_scene_free(vkr_scene_desc* d)
{
    if (!d) return;
    std::free(const_cast<vkr_material*>(d->tri_material));
    std::free(const_cast<vkr_triangle*>(d->tris));
    std::free(const_cast<vkr_sphere*>(d->spheres));
    std::free(const_cast<vkr_plane*>(d->planes));
    std::free(const_cast<vkr_material*>(d->materials));
    *d = vkr_scene_desc{};
}

}  // extern "C"
