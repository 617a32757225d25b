"""GPU parity over a fixed sweep of small random scenes, with the triangle
tree built on the host and on the device: each seed draws its own mix of
sphere materials (diffuse, metal, dielectric, emissive, with random
roughness and metalness), planes, a triangle soup with per-triangle materials
or the reference's single mirror material, an image size (odd sizes
included), samples per pixel, bounce depth and epsilon.  Every pixel must
equal the oracle's linear scan (Assets/Tracer.comp:374-593) bit for bit."""
import numpy as np
import pytest

import oracle_ffi
import scenes
import vkr
from gpu_util import compare, gpu_trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    c = vkr.Context()
    yield c
    c.close()


def _material(rng):
    kind = rng.integers(0, 4)
    a = rng.random(3).astype(np.float32)
    if kind == 0:
        return vkr.material(a, (0, 0, 0), float(rng.uniform(0.05, 1.0)), float(rng.uniform(0.0, 1.0)))
    if kind == 1:
        return vkr.material(a, (0, 0, 0), float(rng.uniform(0.0, 0.3)), 1.0)
    if kind == 2:
        return vkr.material(a, (0, 0, 0), float(rng.uniform(0.3, 0.9)), 0.0, vkr.DIELECTRIC)
    return vkr.material(a, tuple(float(x) for x in rng.uniform(1.0, 64.0) * a), 0.6, 0.0)


def _scene(seed, size=None):
    """size=(w, h) replaces the drawn image size (after every draw, so the
    scene, camera and trace settings stay those of the seed); spp then stays
    at most 2."""
    rng = np.random.default_rng(1000 + seed)
    n_sph = int(rng.integers(2, 40))
    mats = [_material(rng) for _ in range(n_sph)]
    sph = scenes.spheres_from(mats, rng.uniform(-12, 12, (n_sph, 3)).astype(np.float32),
                              rng.uniform(0.3, 3.0, n_sph).astype(np.float32))
    n_pl = int(rng.integers(0, 4))
    normals = [(0, 1, 0), (1, 0, 0), (0, 0, -1), (-1, 0, 0)][:n_pl]
    pln = scenes.planes_from([_material(rng) for _ in range(n_pl)], normals,
                             rng.uniform(12.0, 20.0, n_pl).astype(np.float32))
    n_tri = int(rng.integers(0, 600))
    tris = scenes.random_soup(rng, n_tri, lo=-14.0, hi=14.0, ehi=4.0) if n_tri else np.zeros(0, vkr.TRIANGLE)
    flags, table = 0, np.zeros(0, vkr.MATERIAL)
    if n_tri and rng.random() < 0.5:
        table = np.array([_material(rng) for _ in range(5)], vkr.MATERIAL)
        tris["mat_index"] = rng.integers(0, 5, n_tri).astype(np.uint32)
        flags = vkr.SCENE_TRI_MATERIAL_INDEX
    sc = vkr.Scene(tris=tris, spheres=sph, planes=pln, materials=table, flags=flags)
    pos = rng.uniform(-6, 6, 3) + np.array([0.0, 0.0, -30.0])
    cam = vkr.camera_from_angles(tuple(float(x) for x in pos), float(rng.uniform(-20, 20)),
                                 float(rng.uniform(70, 110)))
    w, h = int(rng.integers(17, 200)), int(rng.integers(13, 150))
    fseed = float(rng.integers(0, 1000))
    spp, depth = int(rng.integers(1, 4)), int(rng.integers(1, 7))
    eps = float(rng.choice([1e-3, 1e-3, 5e-4, 2e-3]))
    if size is not None:
        w, h = size
        spp = min(spp, 2)
    fr = scenes.frame(cam, np.float32(w) / np.float32(h), seed=fseed)
    return scenes.Workload(f"sweep{seed}", sc, fr, w, h, spp, depth), eps


@pytest.mark.parametrize("variant", ["auto", "other"], ids=["kernel_auto", "kernel_other"])
@pytest.mark.parametrize("device_bvh", [False, True], ids=["host_bvh", "device_bvh"])
@pytest.mark.parametrize("seed", range(48))
def test_random_scene_matches_oracle(ctx, torch_cuda, seed, device_bvh, variant):
    """variant "other" forces the kernel variant the scene would not get:
    the wave-shared ray pool below two emitters, single rays from two on."""
    wl, eps = _scene(seed)
    print(f"\n[sweep {seed}] {len(wl.scene.tris)} tris {len(wl.scene.spheres)} spheres {len(wl.scene.planes)} planes "
          f"{wl.width}x{wl.height} spp {wl.spp} depth {wl.max_depth} eps {eps} flags {wl.scene.flags}", flush=True)
    sid = ctx.upload(wl.scene, device_bvh=device_bvh)
    if device_bvh and len(wl.scene.tris) >= 2:
        assert ctx.info(sid)["builder"] == vkr.BUILDER_DEVICE_LBVH
    flags = 0
    if variant == "other":
        flags = vkr.TRACE_RAY_POOL if ctx.info(sid)["n_emitters"] < 2 else vkr.TRACE_NO_RAY_POOL
    g = gpu_trace_eps(ctx, sid, wl, eps, flags)
    r = oracle_ffi.trace(wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth, epsilon=eps)
    st = compare(g, r)
    print(f"[sweep {seed}] " + " ".join(f"{k}={v}" for k, v in st.items()), flush=True)
    assert st["window_overflow"] == 0, st
    assert st["prim_id"] == 1.0 and st["hit_t_bits"] == 1.0 and st["ray_count"] == 1.0, st
    assert st["radiance_1e-4"] == 1.0 and st["rgba8"] == 1.0, st
    ctx.destroy_scene(sid)


@pytest.mark.parametrize("variant", ["auto", "other"], ids=["kernel_auto", "kernel_other"])
@pytest.mark.parametrize("seed", range(6))
def test_random_scene_large_frame(ctx, torch_cuda, seed, variant):
    """The sweep's scenes at 1280x720: an image of more than twice the pixels
    the 4-wave kernel keeps resident runs the 5-wave kernel (24-entry stack),
    which the small frames above do not reach.  4096 random pixels and the
    corners must equal the oracle's linear scan."""
    wl, eps = _scene(seed, size=(1280, 720))
    sid = ctx.upload(wl.scene)
    info = ctx.info(sid)
    print(f"\n[sweep720 {seed}] {len(wl.scene.tris)} tris {len(wl.scene.spheres)} spheres "
          f"spp {wl.spp} depth {wl.max_depth} eps {eps} bvh depth {info['depth']} stack {info['stack_needed']}", flush=True)
    flags = 0
    if variant == "other":
        flags = vkr.TRACE_RAY_POOL if info["n_emitters"] < 2 else vkr.TRACE_NO_RAY_POOL
    g = gpu_trace_eps(ctx, sid, wl, eps, flags)
    W, H = wl.width, wl.height
    pix = np.random.default_rng(300 + seed).choice(W * H, 4096, replace=False).astype(np.uint32)
    pix = np.unique(np.concatenate([pix, np.array([0, W - 1, (H - 1) * W, W * H - 1], np.uint32)]))
    r = oracle_ffi.trace(wl.scene, wl.frame, W, H, wl.spp, wl.max_depth, epsilon=eps, pixels=pix)
    st = compare(g, r, pixels=pix)
    print(f"[sweep720 {seed}] " + " ".join(f"{k}={v}" for k, v in st.items()), flush=True)
    assert st["prim_id"] == 1.0 and st["hit_t_bits"] == 1.0 and st["ray_count"] == 1.0, st
    assert st["radiance_1e-4"] == 1.0 and st["rgba8"] == 1.0, st
    ctx.destroy_scene(sid)


def gpu_trace_eps(ctx, sid, wl, eps, flags=0):
    """gpu_trace with the workload's epsilon set in the trace parameters."""
    orig = wl.params

    def params(**kw):
        p = orig(**kw)
        p.epsilon = eps
        return p

    wl2 = scenes.Workload(wl.name, wl.scene, wl.frame, wl.width, wl.height, wl.spp, wl.max_depth)
    object.__setattr__(wl2, "params", params)
    return gpu_trace(ctx, sid, wl2, flags=flags)


@pytest.mark.parametrize("seed", range(12))
def test_whitted_random_views_match_oracle(ctx, torch_cuda, seed):
    """Raytracer.comp (its own constant scene, Assets/Raytracer.comp:224-355)
    from random cameras, light positions (the push constant it reads) and
    image sizes: bit-exact against the oracle."""
    rng = np.random.default_rng(2000 + seed)
    cam = vkr.camera_from_angles(tuple(float(x) for x in rng.uniform(-40, 40, 3)), float(rng.uniform(-60, 60)),
                                 float(rng.uniform(0, 360)))
    w, h = int(rng.integers(16, 160)), int(rng.integers(16, 160))
    fr = scenes.frame(cam, np.float32(w) / np.float32(h), seed=float(rng.integers(0, 100)),
                      light_pos=tuple(float(x) for x in rng.uniform(-60, 60, 3)))
    wl = scenes.Workload(f"whitted{seed}", scenes.reference_scene(w, h).scene, fr, w, h, 1, 1)
    g = gpu_trace(ctx, 0, wl, kernel=vkr.KERNEL_WHITTED)
    r = oracle_ffi.trace(wl.scene, wl.frame, w, h, kernel=oracle_ffi.KERNEL_WHITTED)
    st = compare(g, r)
    print(f"\n[whitted sweep {seed}] {w}x{h} " + " ".join(f"{k}={v}" for k, v in st.items()), flush=True)
    assert st["prim_id"] == 1.0 and st["hit_t_bits"] == 1.0 and st["rgba8"] == 1.0, st
    assert st["radiance_1e-4"] == 1.0, st


def test_random_scenes_in_one_batch(ctx, torch_cuda):
    """The 48 sweep scenes (sizes, spp, depths, epsilons and emitter counts all
    different) traced as one batched persistent launch: every job equals its
    single-frame trace bit for bit."""
    torch = torch_cuda
    keys = ("rgba8", "radiance", "prim_id", "hit_t", "ray_count")

    def bufs(wl):
        n = wl.width * wl.height
        return {"rgba8": torch.zeros((n, 4), dtype=torch.uint8, device="cuda"),
                "radiance": torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
                "prim_id": torch.zeros(n, dtype=torch.int32, device="cuda"),
                "hit_t": torch.zeros(n, dtype=torch.float32, device="cuda"),
                "ray_count": torch.zeros(n, dtype=torch.int32, device="cuda")}

    cases = [_scene(s) for s in range(48)]
    sids = [ctx.upload(wl.scene) for wl, _ in cases]
    single, batched, jobs = [], [], []
    for (wl, eps), sid in zip(cases, sids):
        p = wl.params()
        p.epsilon = eps
        b = bufs(wl)
        ctx.trace(sid, wl.frame, p, **b)
        single.append(b)
        bb = bufs(wl)
        batched.append(bb)
        jobs.append(ctx.job(sid, wl.frame, p, **bb))
    ctx.trace_batch(jobs)
    torch.cuda.synchronize()
    for i, (s, b) in enumerate(zip(single, batched)):
        for k in keys:
            # Bit patterns: the reference's shading yields NaN on a few
            # pixels (equal bits, but NaN != NaN).
            x, y = s[k], b[k]
            if x.dtype == torch.float32:
                x, y = x.view(torch.int32), y.view(torch.int32)
            assert torch.equal(x, y), (i, k)
    for sid in sids:
        ctx.destroy_scene(sid)


@pytest.mark.parametrize("seed", range(10))
def test_filter_random_sizes_match_oracle(ctx, torch_cuda, seed):
    """Fullscreen.frag as a kernel (bilinear fetch with a black border, the
    temporal-variance gate between the current and the previous traced image,
    v flip) on random traced and present sizes, up- and down-scaling, with
    partly changed previous images: bit-exact against the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(3000 + seed)
    sw, sh = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    dw, dh = int(rng.integers(1, 400)), int(rng.integers(1, 400))
    cur = rng.integers(0, 256, (sw * sh, 4), dtype=np.uint8)
    prev = cur.copy()
    changed = rng.random(sw * sh) < rng.uniform(0.0, 1.0)
    prev[changed] = np.clip(prev[changed].astype(np.int32) + rng.integers(-3, 4, (int(changed.sum()), 4)), 0,
                            255).astype(np.uint8)
    dev = torch.device("cuda:0")
    tc, tp = torch.from_numpy(cur).to(dev), torch.from_numpy(prev).to(dev)
    to = torch.zeros((dw * dh, 4), dtype=torch.uint8, device=dev)
    ctx.filter(tc, tp, sw, sh, to, dw, dh)
    torch.cuda.synchronize()
    ref = np.zeros((dw * dh, 4), np.uint8)
    assert oracle_ffi.load().vkro_filter(cur.ctypes.data, prev.ctypes.data, sw, sh, ref.ctypes.data, dw, dh) == 0
    got = to.cpu().numpy()
    print(f"\n[filter sweep {seed}] {sw}x{sh} -> {dw}x{dh}: {(got == ref).all(axis=1).mean():.6f} equal", flush=True)
    assert np.array_equal(got, ref)
