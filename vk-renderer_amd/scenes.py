"""Seeded scene generators for the benchmark and parity workloads.

The reference renders one fixed scene (4 spheres, 5 planes, 1 triangle; the
constants of Assets/Tracer.comp:186-211 and Source/GraphicsDevice.cpp:817);
`reference_scene` reproduces it.  The other generators build the synthetic
workloads from the reference's own vocabulary (materials per
Assets/Tracer.comp:186-194: "metal" = DIFFUSE with metalness 1 like `mirror`,
"dielectric" = type 1 with the IOR in roughness like `glass`, "emissive" =
DIFFUSE with emissive 128 * albedo like `light`).

Randomness comes from numpy's PCG64 generator seeded with (workload, seed),
so every scene is reproducible bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import vkr

MATTE = dict(roughness=0.4, metalness=0.0)   # matte_* (Tracer.comp:187-189)
LIGHT_E = 128.0                              # light (Tracer.comp:194)


@dataclass
class Workload:
    name: str
    scene: vkr.Scene
    frame: np.ndarray
    width: int
    height: int
    spp: int
    max_depth: int

    def params(self, **kw) -> vkr.TraceParams:
        return vkr.Context.params(self.width, self.height, self.spp, self.max_depth, **kw)


def rng_for(workload_id: int, seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([workload_id, seed]))


def frame(camera: np.ndarray, aspect: float, seed: float = 0.0, light_pos=(0.0, 64.0, 0.0)) -> np.ndarray:
    f = np.zeros((), vkr.FRAME)
    f["aspect_ratio"] = aspect
    f["seed"] = seed
    f["light_pos"] = light_pos
    f["camera"] = camera
    return f


def spheres_from(mats: list, centres, radii) -> np.ndarray:
    s = np.zeros(len(mats), vkr.SPHERE)
    for i, m in enumerate(mats):
        s[i]["mat"] = m
    s["P"] = np.asarray(centres, np.float32).reshape(-1, 3)
    s["r"] = np.asarray(radii, np.float32)
    return s


def planes_from(mats: list, normals, lens) -> np.ndarray:
    p = np.zeros(len(mats), vkr.PLANE)
    for i, m in enumerate(mats):
        p[i]["mat"] = m
    p["N"] = np.asarray(normals, np.float32).reshape(-1, 3)
    p["len"] = np.asarray(lens, np.float32)
    return p


def tris_from(v0, v1, v2, mat_index=None) -> np.ndarray:
    t = np.zeros(len(v0), vkr.TRIANGLE)
    t["v0"] = v0
    t["v1"] = v1
    t["v2"] = v2
    if mat_index is not None:
        t["mat_index"] = mat_index
    return t


# --------------------------------------------------------------------------
# The reference scene (Assets/Tracer.comp:186-211, GraphicsDevice.cpp:817,
# camera of Source/Main.cpp:134-141, aspect of GraphicsDevice.cpp:1260 with the
# 1024x768 window of Main.cpp:92, RES 1024 of Main.cpp:112)
# --------------------------------------------------------------------------
def reference_materials() -> dict:
    m = vkr.material
    return {
        "matte_white": m((1.0, 1.0, 1.0), (0, 0, 0), 0.3, 0.7),
        "matte_red": m((0.75, 0.25, 0.25), (0, 0, 0), 0.4, 0.0),
        "matte_green": m((0.25, 0.75, 0.25), (0, 0, 0), 0.4, 0.0),
        "matte_blue": m((0.25, 0.25, 0.75), (0, 0, 0), 0.4, 0.0),
        "plastic": m((0.25, 0.25, 0.75), (0, 0, 0), 0.3, 0.6),
        "mirror": m((1.0, 0.5, 0.5), (0, 0, 0), 0.0, 1.0),
        "glass": m((1.0, 1.0, 1.0), (0, 0, 0), 0.42, 0.0, vkr.DIELECTRIC),
        "light": m((1.0, 1.0, 1.0), (128.0, 128.0, 128.0), 0.6, 0.0),
    }


def reference_scene(width: int = 1024, height: int = 1024, spp: int = 4, max_depth: int = 4,
                    seed: float = 0.0) -> Workload:
    M = reference_materials()
    spheres = spheres_from([M["glass"], M["light"], M["mirror"], M["plastic"]],
                           [(42.0, 16.0, 12.0), (0.0, 96.0, 0.0), (-32.0, 24.0, 24.0), (-24.0, 11.0, -48.0)],
                           [16.0, 12.0, 24.0, 11.0])
    planes = planes_from([M["matte_white"], M["matte_white"], M["matte_red"], M["matte_green"], M["matte_blue"]],
                         [(0, 1, 0), (0, -1, 0), (1, 0, 0), (0, 0, -1), (-1, 0, 0)],
                         [0.0, 128.0, 64.0, 64.0, 64.0])
    tris = tris_from([(10.0, 10.0, 0.0)], [(0.0, 20.0, 0.0)], [(-10.0, 10.0, 0.0)])
    cam = vkr.camera_from_angles((32.8509, 30.6991, -106.389), 4.44998, -602.79)
    f = frame(cam, np.float32(1024.0) / np.float32(768.0), seed)
    sc = vkr.Scene(tris=tris, spheres=spheres, planes=planes, tri_material=M["mirror"])
    return Workload("reference_scene", sc, f, width, height, spp, max_depth)


# --------------------------------------------------------------------------
# Random content
# --------------------------------------------------------------------------
def random_sphere_materials(rng: np.random.Generator, n: int, kinds=("diffuse", "metal", "dielectric", "emissive")):
    kind = rng.integers(0, len(kinds), n)
    albedo = rng.random((n, 3), dtype=np.float64).astype(np.float32)
    mats = []
    for i in range(n):
        k = kinds[kind[i]]
        a = albedo[i]
        if k == "diffuse":
            mats.append(vkr.material(a, (0, 0, 0), **MATTE))
        elif k == "metal":
            mats.append(vkr.material(a, (0, 0, 0), 0.0, 1.0))
        elif k == "dielectric":
            mats.append(vkr.material(a, (0, 0, 0), 0.42, 0.0, vkr.DIELECTRIC))
        else:
            mats.append(vkr.material(a, LIGHT_E * a, 0.6, 0.0))
    return mats


def random_spheres(rng, n, lo, hi, rlo=0.3, rhi=1.5, kinds=("diffuse", "metal", "dielectric", "emissive"),
                   ylo=None, yhi=None):
    mats = random_sphere_materials(rng, n, kinds)
    c = rng.uniform(lo, hi, (n, 3))
    if ylo is not None:
        c[:, 1] = rng.uniform(ylo, yhi, n)
    r = rng.uniform(rlo, rhi, n)
    return spheres_from(mats, c.astype(np.float32), r.astype(np.float32))


def light_sphere(centre, radius, e=LIGHT_E) -> np.ndarray:
    return spheres_from([vkr.material((1.0, 1.0, 1.0), (e, e, e), 0.6, 0.0)], [centre], [radius])


def ground(y: float) -> np.ndarray:
    M = reference_materials()
    return planes_from([M["matte_white"]], [(0.0, 1.0, 0.0)], [-y])


def random_soup(rng, n, lo=-50.0, hi=50.0, elo=0.1, ehi=2.0) -> np.ndarray:
    v0 = rng.uniform(lo, hi, (n, 3))

    def unit(k):
        u = rng.normal(size=(k, 3))
        return u / np.linalg.norm(u, axis=1, keepdims=True)

    v1 = v0 + rng.uniform(elo, ehi, (n, 1)) * unit(n)
    v2 = v0 + rng.uniform(elo, ehi, (n, 1)) * unit(n)
    return tris_from(v0.astype(np.float32), v1.astype(np.float32), v2.astype(np.float32))


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------
def spheres64(seed: int = 0) -> Workload:
    """512x512, 1 spp, 3 bounces, 64 spheres in [-10,10]^3 + ground plane."""
    rng = rng_for(1, seed)
    sph = random_spheres(rng, 64, -10.0, 10.0)
    cam = vkr.camera_from_angles((0.0, 0.0, -30.0), 0.0, 90.0)
    sc = vkr.Scene(spheres=sph, planes=ground(-11.0))
    return Workload("spheres64_512", sc, frame(cam, 1.0), 512, 512, 1, 3)


def soup100k(seed: int = 0, n_tris: int = 100_000, width: int = 1920, height: int = 1080, spp: int = 4,
             max_depth: int = 4, n_spheres: int = 256) -> Workload:
    """1920x1080, 4 spp, 4 bounces, 100k random triangles + 256 spheres + 1 light."""
    rng = rng_for(2, seed)
    tris = random_soup(rng, n_tris)
    sph = np.concatenate([random_spheres(rng, n_spheres, -10.0, 10.0), light_sphere((0.0, 80.0, 0.0), 12.0)])
    cam = vkr.camera_from_angles((0.0, 0.0, -120.0), 0.0, 90.0)
    sc = vkr.Scene(tris=tris, spheres=sph, planes=ground(-55.0))
    return Workload(f"soup{n_tris}_{width}x{height}", sc, frame(cam, np.float32(width) / np.float32(height)),
                    width, height, spp, max_depth)


def many_emitters(n_spheres: int = 140, seed: int = 0, width: int = 64, height: int = 48, spp: int = 2,
                  max_depth: int = 3) -> Workload:
    """Spheres of which about half emit (diffuse / metal / emissive /
    emissive), a 400-triangle soup and a ground plane: with more than 32
    emitters a diffuse hit's next-event samples span several evaluation
    windows of the pooled kernel (seed 0: 69 emitters; n_spheres 66, seed 3:
    33; n_spheres 60, seed 5: 32)."""
    rng = rng_for(11, seed)
    sph = random_spheres(rng, n_spheres, -12.0, 12.0, kinds=("diffuse", "metal", "emissive", "emissive"))
    tris = random_soup(rng, 400, lo=-15.0, hi=15.0)
    cam = vkr.camera_from_angles((0.0, 0.0, -40.0), 0.0, 90.0)
    sc = vkr.Scene(tris=tris, spheres=sph, planes=ground(-13.0))
    return Workload(f"many_emitters_{n_spheres}_{seed}", sc, frame(cam, np.float32(width) / np.float32(height)),
                    width, height, spp, max_depth)


def far_sphere(res: int = 128, dist: float = 2000.0, span: float = 1.5e-3) -> Workload:
    """A unit sphere seen from `dist` away through a narrow camera (the image
    spans +-span*dist around it), plus a mirror-triangle soup and two more
    spheres off screen.  From that far the reference's sphere test
    (Assets/Tracer.comp:314-330) rounds its discriminant h = b^2 - 4c with an
    error of tens of ulp of |o - c|^2, so it reports hits on rays that pass
    up to ~0.2 outside the sphere: a traversal that culls by the exact box
    loses them."""
    rng = rng_for(7, 0)
    M = reference_materials()
    sph = spheres_from([M["matte_white"], M["plastic"], M["light"]],
                       [(0.3, -0.2, 0.1), (40.0, 5.0, 30.0), (0.0, 300.0, -1500.0)], [1.0, 2.0, 40.0])
    tris = random_soup(rng, 64, lo=-60.0, hi=-20.0)
    cam = np.zeros((), vkr.CAMERA)
    cam["pos"] = (0.0, 0.0, -dist)
    cam["dir"] = (0.0, 0.0, 1.0)
    cam["right"] = (span, 0.0, 0.0)
    cam["up"] = (0.0, span, 0.0)
    sc = vkr.Scene(tris=tris, spheres=sph, tri_material=M["mirror"])
    return Workload(f"far_sphere_{res}", sc, frame(cam, 1.0), res, res, 1, 2)


def value_noise_fbm(rng, x, z, octaves=6, lacunarity=2.0, gain=0.5, base_freq=1.0 / 128.0, lattice=512):
    table = rng.uniform(-1.0, 1.0, (lattice, lattice))
    total = np.zeros_like(x)
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        a = table[ix % lattice, iz % lattice]
        b = table[(ix + 1) % lattice, iz % lattice]
        c = table[ix % lattice, (iz + 1) % lattice]
        d = table[(ix + 1) % lattice, (iz + 1) % lattice]
        total += amp * ((a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def heightfield(seed: int = 0, n: int = 1024, extent: float = 512.0, amplitude: float = 20.0, n_spheres: int = 4096,
                width: int = 1920, height: int = 1080, spp: int = 8, max_depth: int = 8) -> Workload:
    """1920x1080, 8 spp, 8 bounces: 1024^2-vertex fBm heightfield + 4096 metal/dielectric spheres + 1 light."""
    rng = rng_for(3, seed)
    xs = np.linspace(-extent, extent, n)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = amplitude * value_noise_fbm(rng, X + extent, Z + extent)
    V = np.stack([X, Y, Z], axis=-1).astype(np.float32)
    a, b = V[:-1, :-1].reshape(-1, 3), V[1:, :-1].reshape(-1, 3)      # (i,j), (i+1,j)
    c, d = V[:-1, 1:].reshape(-1, 3), V[1:, 1:].reshape(-1, 3)        # (i,j+1), (i+1,j+1)
    # Counter-clockwise seen from +y, so the one-sided test accepts rays from above.
    v0 = np.concatenate([a, b])
    v1 = np.concatenate([c, c])
    v2 = np.concatenate([b, d])
    tris = tris_from(v0, v1, v2)
    sph = random_spheres(rng, n_spheres, -extent, extent, kinds=("metal", "dielectric"), ylo=25.0, yhi=60.0)
    sph = np.concatenate([sph, light_sphere((0.0, 300.0, 0.0), 60.0)])
    cam = vkr.camera_from_angles((0.0, 90.0, -600.0), -10.0, 90.0)
    sc = vkr.Scene(tris=tris, spheres=sph)
    return Workload(f"heightfield{len(tris)}_{width}x{height}", sc, frame(cam, np.float32(width) / np.float32(height)),
                    width, height, spp, max_depth)


def adjoint_scene(seed: int = 0, width: int = 256, height: int = 256, spp: int = 16, max_depth: int = 4) -> Workload:
    """256x256, 16 spp, 4 bounces: 32 diffuse/metal spheres + 1 light sphere (+ ground)."""
    rng = rng_for(5, seed)
    sph = random_spheres(rng, 32, -10.0, 10.0, kinds=("diffuse", "metal"))
    sph = np.concatenate([sph, light_sphere((0.0, 30.0, 0.0), 6.0)])
    cam = vkr.camera_from_angles((0.0, 0.0, -30.0), 0.0, 90.0)
    sc = vkr.Scene(spheres=sph, planes=ground(-11.0))
    return Workload("adjoint32", sc, frame(cam, np.float32(width) / np.float32(height)), width, height, spp, max_depth)


BATCH_RES = ((256, 256), (512, 512), (1280, 720), (1920, 1080))


def batch_jobs(n_scenes: int = 64, seed: int = 0, max_tris: int = 1_000_000) -> list[Workload]:
    """64 scenes: triangle counts log-uniform in [1e3, 1e6], resolutions from BATCH_RES, 1..8 spp, 4 bounces."""
    rng = rng_for(4, seed)
    out = []
    for i in range(n_scenes):
        n = int(round(10.0 ** rng.uniform(3.0, np.log10(max_tris))))
        w, h = BATCH_RES[int(rng.integers(0, len(BATCH_RES)))]
        spp = int(rng.integers(1, 9))
        wl = soup100k(seed=1000 + i, n_tris=n, width=w, height=h, spp=spp, max_depth=4)
        wl.frame["seed"] = np.float32(0.0)
        out.append(wl)
    return out


def sphere_params(spheres: np.ndarray, light_scale: float = 1.0) -> np.ndarray:
    """Packs sphere materials into the 7 * n + 1 parameter vector of vkr_scene_set_sphere_params."""
    n = len(spheres)
    p = np.zeros(7 * n + 1, np.float32)
    for i in range(n):
        m = spheres[i]["mat"]
        p[7 * i:7 * i + 3] = m["albedo"]
        p[7 * i + 3:7 * i + 6] = m["emissive"] / np.float32(light_scale)
        p[7 * i + 6] = m["roughness"]
    p[7 * n] = light_scale
    return p
