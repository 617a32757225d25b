"""GPU: vkr_trace_frame_static, the entry point that allocates nothing (its
kernels' scratch is a static device array), against vkr_trace_frame on the
same inputs: bit-identical outputs and counters on two calls in a row, the
second with a smaller image, for a scene traced with the wave-shared ray
pool (per-lane records in the static array, grid capped to them) and one
traced one ray per lane."""
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


def _bufs(torch, n):
    dev = torch.device("cuda:0")
    return {"rgba8": torch.zeros((n, 4), dtype=torch.uint8, device=dev),
            "radiance": torch.zeros((n, 4), dtype=torch.float32, device=dev),
            "prim_id": torch.zeros(n, dtype=torch.int32, device=dev),
            "hit_t": torch.zeros(n, dtype=torch.float32, device=dev),
            "ray_count": torch.zeros(n, dtype=torch.int32, device=dev)}


def _same(torch, x, y):
    if x.dtype == torch.float32:
        x, y = x.view(torch.int32), y.view(torch.int32)
    return torch.equal(x, y)


@pytest.mark.parametrize("kind", ["ray_pool", "single_ray"])
def test_static_entry_matches_trace_frame(torch_cuda, kind):
    torch = torch_cuda
    if kind == "ray_pool":     # 58 emitters: per-lane records in the static array
        wl = scenes.soup100k(n_tris=20000, width=640, height=360, spp=2)
    else:                      # 1 emitter: one ray per lane, no records
        wl = scenes.reference_scene(512, 384, spp=2, max_depth=4)
    with vkr.Context() as ctx:
        sid = ctx.upload(wl.scene)
        assert (ctx.info(sid)["n_emitters"] >= 2) == (kind == "ray_pool")
        # Two calls in a row, the second on a smaller image: the static
        # array keeps the first call's contents, which must not leak.
        for w, h in ((wl.width, wl.height), (wl.width // 2 - 3, wl.height // 2 + 5)):
            p = vkr.Context.params(w, h, wl.spp, wl.max_depth)
            ref = _bufs(torch, w * h)
            ctx.trace(sid, wl.frame, p, **ref)
            torch.cuda.synchronize()
            c_ref = ctx.counters()
            got = _bufs(torch, w * h)
            ctx.trace_static(sid, wl.frame, p, **got)
            c_got = ctx.counters()
            print(f"\n[static {kind} {w}x{h}] rays {c_got['rays']} traced {c_got['rays_traced']}", flush=True)
            for k in KEYS:
                assert _same(torch, ref[k], got[k]), (kind, w, h, k)
            for k in ("rays", "shadow_rays", "rays_traced", "window_overflow"):
                assert c_ref[k] == c_got[k], (kind, w, h, k)
        # A one-pixel-per-lane launch with the ray pool needs a record per
        # pixel: refused when they do not fit the static array.
        if kind == "ray_pool":
            big = vkr.Context.params(4096, 4096, 1, 1, sched=vkr.SCHED_PLAIN)
            with pytest.raises(vkr.VkrError) as err:
                ctx.trace_static(sid, wl.frame, big)
            assert err.value.code == 4
        ctx.destroy_scene(sid)
