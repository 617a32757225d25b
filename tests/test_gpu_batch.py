"""GPU: batched persistent launches and captured graphs reproduce the
single-frame results bit for bit; the autotuner returns a valid shape."""
import numpy as np
import pytest

import scenes
import vkr

pytestmark = pytest.mark.gpu

KEYS = ("rgba8", "radiance", "prim_id", "hit_t", "ray_count")


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


def _same(torch, x, y):
    """Bitwise equality (float outputs compared as bit patterns: a NaN pixel
    of the reference's shading equals itself)."""
    if x.dtype == torch.float32:
        x, y = x.view(torch.int32), y.view(torch.int32)
    return torch.equal(x, y)


def _bufs(torch, wl):
    dev = torch.device("cuda:0")
    n = wl.width * wl.height
    return {"rgba8": torch.zeros((n, 4), dtype=torch.uint8, device=dev),
            "radiance": torch.zeros((n, 4), dtype=torch.float32, device=dev),
            "prim_id": torch.zeros(n, dtype=torch.int32, device=dev),
            "hit_t": torch.zeros(n, dtype=torch.float32, device=dev),
            "ray_count": torch.zeros(n, dtype=torch.int32, device=dev)}


def _small_jobs():
    out = []
    for i, (n, w, h, spp) in enumerate(((1500, 64, 48, 1), (8000, 96, 96, 3), (300, 40, 72, 2), (20000, 128, 64, 2))):
        wl = scenes.soup100k(seed=50 + i, n_tris=n, width=w, height=h, spp=spp, max_depth=4, n_spheres=24)
        out.append(wl)
    out.append(scenes.reference_scene(80, 60, spp=2, max_depth=4))
    return out


def test_batch_matches_single_frames(ctx, torch_cuda):
    torch = torch_cuda
    wls = _small_jobs()
    sids = [ctx.upload(wl.scene) for wl in wls]
    single = []
    for wl, sid in zip(wls, sids):
        b = _bufs(torch, wl)
        ctx.trace(sid, wl.frame, wl.params(), **b)
        single.append(b)
    torch.cuda.synchronize()
    single_rays = 0
    for wl, sid in zip(wls, sids):
        ctx.trace(sid, wl.frame, wl.params(), **_bufs(torch, wl))
        single_rays += ctx.counters()["rays"]
    for block, chunk in ((128, 32), (64, 1), (256, 64)):
        batched = [_bufs(torch, wl) for wl in wls]
        jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, batched)]
        ctx.trace_batch(jobs, block=block, chunk=chunk)
        torch.cuda.synchronize()
        assert ctx.counters()["rays"] == single_rays
        for s, b in zip(single, batched):
            for k in KEYS:
                assert _same(torch, s[k], b[k]), (block, chunk, k)


def test_batch_mixed_jobs_in_waves(ctx, torch_cuda):
    """Eight multi-emitter scenes of very different sizes (pooled kernel), small
    images and a refill threshold of 1: waves hold pixels of several jobs and
    trace each other's shadow rays, each against its own job's scene."""
    torch = torch_cuda
    wls = []
    for i, n in enumerate((200, 40000, 1000, 12000, 300, 60000, 5000, 800)):
        wl = scenes.soup100k(seed=80 + i, n_tris=n, width=24 + 8 * (i % 3), height=16 + 8 * (i % 2), spp=2,
                             max_depth=4, n_spheres=24)
        wls.append(wl)
    sids = [ctx.upload(wl.scene) for wl in wls]
    assert all(ctx.info(sid)["n_emitters"] >= 2 for sid in sids)
    single = []
    for wl, sid in zip(wls, sids):
        b = _bufs(torch, wl)
        ctx.trace(sid, wl.frame, wl.params(), **b)
        single.append(b)
    torch.cuda.synchronize()
    for block, chunk in ((64, 1), (128, 1), (256, 4)):
        batched = [_bufs(torch, wl) for wl in wls]
        jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, batched)]
        ctx.trace_batch(jobs, block=block, chunk=chunk)
        torch.cuda.synchronize()
        for s, b in zip(single, batched):
            for k in KEYS:
                assert _same(torch, s[k], b[k]), (block, chunk, k)


def test_batch_device_built_scenes(ctx, torch_cuda):
    """A batch mixing scenes whose triangle trees were built on the device
    (VKR_SCENE_DEVICE_BVH) with host-built ones: every job's image equals
    the single-frame trace of its host-built scene."""
    torch = torch_cuda
    wls = [scenes.soup100k(seed=90 + i, n_tris=n, width=24 + 8 * (i % 3), height=16 + 8 * (i % 2), spp=2,
                           max_depth=4, n_spheres=24)
           for i, n in enumerate((3000, 40000, 500, 20000, 7000, 1200))]
    host = [ctx.upload(wl.scene) for wl in wls]
    mixed = [ctx.upload(wl.scene, device_bvh=(i % 2 == 0)) for i, wl in enumerate(wls)]
    assert [ctx.info(s)["builder"] for s in mixed] == [vkr.BUILDER_DEVICE_LBVH, vkr.BUILDER_HOST_SAH] * 3
    single = []
    for wl, sid in zip(wls, host):
        b = _bufs(torch, wl)
        ctx.trace(sid, wl.frame, wl.params(), **b)
        single.append(b)
    batched = [_bufs(torch, wl) for wl in wls]
    jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, mixed, batched)]
    ctx.trace_batch(jobs)
    torch.cuda.synchronize()
    for i, (s, b) in enumerate(zip(single, batched)):
        for k in KEYS:
            assert _same(torch, s[k], b[k]), (i, k)
    for sid in host + mixed:
        ctx.destroy_scene(sid)


def test_graph_replays_match_eager(ctx, torch_cuda):
    torch = torch_cuda
    wls = _small_jobs()[:3]
    sids = [ctx.upload(wl.scene) for wl in wls]
    eager = [_bufs(torch, wl) for wl in wls]
    jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, eager)]
    ctx.trace_batch(jobs)
    torch.cuda.synchronize()
    eager_rays = ctx.counters()["rays"]
    graphed = [_bufs(torch, wl) for wl in wls]
    gjobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, graphed)]
    gid = ctx.batch_graph(gjobs)
    for rep in range(3):
        for b in graphed:
            for v in b.values():
                v.zero_()
        ctx.graph_launch(gid)
        torch.cuda.synchronize()
        assert ctx.counters()["rays"] == eager_rays, rep
        for e, g in zip(eager, graphed):
            for k in KEYS:
                assert _same(torch, e[k], g[k]), (rep, k)
        # unrelated eager work between replays
        ctx.trace(sids[0], wls[0].frame, wls[0].params(), **_bufs(torch, wls[0]))
        torch.cuda.synchronize()
    ctx.graph_destroy(gid)


def test_autotune_picks_a_valid_shape(ctx, torch_cuda):
    torch = torch_cuda
    wl = scenes.reference_scene(128, 96, spp=1, max_depth=4)
    sid = ctx.upload(wl.scene)
    b = _bufs(torch, wl)
    res = ctx.autotune(sid, wl.frame, wl.params(), repeats=1, rgba8=b["rgba8"])
    assert res["block"] in (64, 128, 256) and res["stack"] in (24, 32, 64)
    assert len(res["table"]) >= 10
    assert abs(res["best_ms"] - min(t["ms"] for t in res["table"])) < 1e-3
    # The tuned (AUTO) launch is bit-identical to an explicit plain launch.
    a = _bufs(torch, wl)
    ctx.trace(sid, wl.frame, wl.params(), **a)
    p = _bufs(torch, wl)
    ctx.trace(sid, wl.frame, wl.params(sched=vkr.SCHED_PLAIN, block=64), **p)
    torch.cuda.synchronize()
    for k in KEYS:
        assert _same(torch, a[k], p[k]), k


def test_graph_set_frame_replays_new_push_constants(ctx, torch_cuda):
    """A captured launch replayed with new per-frame data (camera, aspect,
    seed) traces exactly what vkr_trace_frame traces with that data."""
    torch = torch_cuda
    wl = scenes.spheres64()
    sid = ctx.upload(wl.scene)
    b = _bufs(torch, wl)
    gid = ctx.batch_graph([ctx.job(sid, wl.frame, wl.params(), **b)])
    try:
        for k, (pos, pitch, yaw, seed) in enumerate((((0.0, 0.0, -30.0), 0.0, 90.0, 0.25),
                                                     ((4.0, 3.0, -28.0), -8.0, 97.0, 0.75))):
            fr = wl.frame.copy()
            fr["camera"] = vkr.camera_from_angles(pos, pitch, yaw)
            fr["seed"] = np.float32(seed)
            fr["aspect_ratio"] = np.float32(1.25 + k)
            ctx.graph_set_frame(gid, 0, fr)
            ctx.graph_launch(gid)
            ref = _bufs(torch, wl)
            ctx.trace(sid, fr, wl.params(), **ref)
            torch.cuda.synchronize()
            for key in KEYS:
                assert _same(torch, b[key], ref[key]), (k, key)
    finally:
        ctx.graph_destroy(gid)


def test_autotune_batch_is_bit_identical_to_default(ctx, torch_cuda):
    """The batch autotuner (workgroup size, pool chunk, stack depth, chosen
    for this job set) only changes the launch shape: the tuned batch, eager
    and captured, equals the default-shape batch bit for bit."""
    torch = torch_cuda
    wls = _small_jobs()
    sids = [ctx.upload(wl.scene) for wl in wls]
    default = [_bufs(torch, wl) for wl in wls]
    ctx.trace_batch([ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, default)],
                    block=128, chunk=8)
    tuned = [_bufs(torch, wl) for wl in wls]
    jobs = [ctx.job(sid, wl.frame, wl.params(), **b) for wl, sid, b in zip(wls, sids, tuned)]
    res = ctx.autotune_batch(jobs, repeats=1)
    print(f"\n[autotune batch] block {res['block']} chunk {res['chunk']} stack {res['stack']} "
          f"best {res['best_ms']:.3f} ms over {len(res['table'])} candidates", flush=True)
    assert res["block"] in (64, 128, 256) and res["chunk"] in (1, 4, 8, 16, 32) and res["stack"] in (24, 32, 64)
    assert len(res["table"]) >= 7
    assert {t["block"] for t in res["table"]} == {64, 128, 256}
    assert abs(res["best_ms"] - min(t["ms"] for t in res["table"])) < 1e-3
    for b in tuned:
        for v in b.values():
            v.zero_()
    ctx.trace_batch(jobs)                      # block = chunk = 0: the tuned shape
    torch.cuda.synchronize()
    for d, t in zip(default, tuned):
        for k in KEYS:
            assert _same(torch, d[k], t[k]), k
    for b in tuned:
        for v in b.values():
            v.zero_()
    gid = ctx.batch_graph(jobs)
    ctx.graph_launch(gid)
    torch.cuda.synchronize()
    for d, t in zip(default, tuned):
        for k in KEYS:
            assert _same(torch, d[k], t[k]), k
    ctx.graph_destroy(gid)
    for sid in sids:
        ctx.destroy_scene(sid)


def test_graph_follows_sphere_parameter_changes(ctx, torch_cuda):
    """A captured graph keeps tracing correctly when vkr_scene_set_sphere_params
    changes the scene's emitter list (1 -> 3 -> 6 -> 0 emitters): each replay
    equals vkr_trace_frame with the same parameters.  The scene cannot be
    destroyed while the graph exists."""
    torch = torch_cuda
    rng = np.random.default_rng(21)
    sph = scenes.random_spheres(rng, 10, -6.0, 6.0, kinds=("diffuse", "metal"))
    sc = vkr.Scene(spheres=sph, planes=scenes.ground(-7.0))
    theta = scenes.sphere_params(sph, light_scale=1.0)
    n_sph = len(sph)

    def with_emitters(k):
        t = theta.copy()
        for s in range(n_sph):
            t[7 * s + 3:7 * s + 6] = (40.0, 30.0, 20.0) if s < k else (0.0, 0.0, 0.0)
        return t

    cam = vkr.camera_from_angles((0.0, 2.0, -22.0), -5.0, 90.0)
    wl = scenes.Workload("emitter_changes", sc, scenes.frame(cam, np.float32(64) / np.float32(48)), 64, 48, 2, 4)
    sid = ctx.upload(sc)
    ctx.set_sphere_params(sid, with_emitters(1))
    b = _bufs(torch, wl)
    gid = ctx.batch_graph([ctx.job(sid, wl.frame, wl.params(), **b)])
    try:
        with pytest.raises(vkr.VkrError) as err:
            ctx.destroy_scene(sid)
        assert err.value.code == 4
        for k in (1, 3, 6, 0):
            ctx.set_sphere_params(sid, with_emitters(k))
            assert ctx.info(sid)["n_emitters"] == k
            for v in b.values():
                v.zero_()
            ctx.graph_launch(gid)
            ref = _bufs(torch, wl)
            ctx.trace(sid, wl.frame, wl.params(), **ref)
            torch.cuda.synchronize()
            for key in KEYS:
                assert _same(torch, b[key], ref[key]), (k, key)
    finally:
        ctx.graph_destroy(gid)
    ctx.destroy_scene(sid)


def test_batch_pool_cursor_limit(ctx, torch_cuda):
    """The 32-bit pool cursor can run past the pool's end by the resident
    lanes: a batch whose entries come within that margin of 2^32 is refused
    before anything is launched (eager and captured)."""
    wl = scenes.reference_scene(8, 8, spp=1, max_depth=1)
    sid = ctx.upload(wl.scene)
    big = vkr.Context.params(16384, 16384, 1, 1)          # 2^28 pool entries
    last = vkr.Context.params(16384, 16376, 1, 1)         # 2^28 - 2^17
    jobs = [ctx.job(sid, wl.frame, big) for _ in range(15)] + [ctx.job(sid, wl.frame, last)]
    for call in (lambda: ctx.trace_batch(jobs, block=128), lambda: ctx.batch_graph(jobs, block=128)):
        with pytest.raises(vkr.VkrError) as err:
            call()
        assert err.value.code == 4 and "2^32" in str(err.value)
    ctx.destroy_scene(sid)
